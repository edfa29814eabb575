/*
 * snch.h -- C ABI of the MI355X (gfx950) SNCH-LBVH library (libsnch.so).
 *
 * The reference (tyanyuy3125/snch-lbvh) is a header-only CUDA template
 * library; its host entry points are lbvh::scene<2|3> and lbvh::bvh, and its
 * device entry points are query_device()/sample_object_in_sphere() called
 * one thread per query from the user's own kernels.  This ABI replaces both
 * with plain C: host-side scene/build handles and batched query launches on a
 * caller-supplied HIP stream.  No C++ or HIP types cross the boundary.
 *
 * Conventions
 *  - Every function returns SNCH_OK (0) or an SNCH_ERR_* code; the message of
 *    the most recent failure on the calling thread is snch_last_error().
 *  - Query inputs and outputs are DEVICE pointers (hipMalloc'ed or torch CUDA
 *    tensors); mesh inputs to snch_scene_create_* are HOST pointers.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  Query
 *    launches are asynchronous; builds complete before returning.
 *  - A scene handle is externally synchronised: one thread at a time per
 *    handle; different handles may be used from different threads.
 *  - Index sentinels follow the reference: 0xFFFFFFFF for "no primitive"
 *    (query.cuh:254), +inf distance (query.cuh:343), -1 for "no sample"
 *    (sample.cuh:255).
 */
#ifndef SNCH_H
#define SNCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNCH_OK 0
#define SNCH_ERR_INVALID 1  /* bad argument (null pointer, index out of range, NaN vertex) */
#define SNCH_ERR_EMPTY 2    /* empty mesh: the reference silently returns (bvh.cuh:383-386) */
#define SNCH_ERR_HIP 3      /* HIP runtime failure */
#define SNCH_ERR_STATE 4    /* "BVH is not built yet." (scene.cuh:686, :1250) */
#define SNCH_ERR_OVERFLOW 5 /* a traversal exceeded its stack or its step bound, which only a
                               malformed tree can do (query.cuh:315, :420 assert) */

typedef struct snch_context snch_context;
typedef struct snch_scene snch_scene;

/* Build flags */
#define SNCH_BUILD_REFERENCE_RADIUS 1u /* use the reference's cone radius (midpoints, sqrt merge);
                                          default is the enclosing radius, which never prunes a
                                          true nearest silhouette */
#define SNCH_BUILD_TIME_PHASES 2u      /* record GPU events between the build's phases (read them
                                          with snch_scene_build_phases); the tree is unchanged */
/* Layout of the refit above 16384-leaf subtrees (one workgroup): by default
 * the nodes' boxes sit in LDS while they fit, their topology while it fits,
 * else the merge climbs with arrival counters in memory.  These two force the
 * second or third layout at any size (for tests); the tree is unchanged. */
#define SNCH_BUILD_TOPS_IN_MEMORY 4u
#define SNCH_BUILD_TOPS_CLIMB 8u

typedef struct snch_build_opts {
    uint32_t flags;     /* SNCH_BUILD_* */
    uint32_t leaf_size; /* 1 = the reference topology (bvh.cuh:200-229); 0 means 1 */
} snch_build_opts;

#define SNCH_ORDER_MIN_QUERIES 65536u

/* Traversal tuning (chosen by the autotuner; defaults are safe) */
typedef struct snch_query_opts {
    uint32_t block_size; /* threads per block: 64, 128, 256 or 512; 0 = default (64).  The
                            closest silhouette without a seed (or with a hard cap) runs, at
                            block size 64 and stack depth 16, as a persistent launch whose
                            wavefronts take 64-query chunks; other shapes launch one block
                            per chunk.  Outputs are the same. */
    uint32_t coherent;   /* visiting order: 0 = Morton order of the query points when
                            n >= SNCH_ORDER_MIN_QUERIES, 1 = always Morton order,
                            2 = input order.  Results are written in input order and
                            do not depend on the visiting order. */
    uint64_t *stats;     /* optional DEVICE uint64[5], accumulated: child box tests, leaf
                            evaluations, primitive / silhouette tests, internal nodes
                            visited, cone records read (silhouette query only) */
    const uint32_t *order; /* optional DEVICE permutation of [0, n) (e.g. from
                              snch_query_order): the visiting order; overrides coherent */
    uint32_t stack_depth;      /* LDS stack entries per lane: closest point 8 or 16 (0 = 16),
                                  closest silhouette (3D) and batched queries 8, 16, 24 or
                                  32 (0 = 16) */
    uint32_t queries_per_wave; /* batched queries: queries a wavefront takes from the work
                                  queue at a time, a multiple of 64 up to 4096 (0 = 64) */
    uint32_t seed_is_cap;      /* closest silhouette: 1 = seed_radius is a hard search radius
                                  (the answer is exact when it lies within the radius; otherwise
                                  the returned distance exceeds the radius, possibly +inf) */
    uint32_t seed_retry;       /* closest silhouette with seed_radius (not a cap): 0 = queries not
                                  settled within their radius are queued (one queue per XCD) and
                                  searched again 64 at a time by whole wavefronts of the same
                                  persistent launch (default;
                                  block size 64 and stack depth 16 only), 1 = each is searched again
                                  at once by its own lane.  Outputs are the same. */
} snch_query_opts;

const char *snch_last_error(void);
int snch_version(void);

/* ---- context ------------------------------------------------------------ */
int snch_context_create(int device, snch_context **out);
int snch_context_destroy(snch_context *ctx);

/* ---- scenes ------------------------------------------------------------- */
/* scene<3>(vertices, indices) + compute_silhouettes() (scene.cuh:1128-1204).
 * V: nV x 3 floats, F: nF x 3 int32 (host).  Builds the edge table with the
 * reference's edge ids (first appearance over triangle i, local edge
 * (v_j, v_{j+1})), edge slots (opp0, lo, hi, opp1) and edge ownership.
 * At most 2^24 triangles (segments for scene<2>). */
int snch_scene_create_tri(snch_context *ctx, const float *V, uint32_t nV, const int32_t *F, uint32_t nF,
                          snch_scene **out);
/* scene<2>(vertices, indices) + compute_silhouettes() (scene.cuh:624-652).
 * V: nV x 2 floats, S: nS x 2 int32 line segments (host). */
int snch_scene_create_seg(snch_context *ctx, const float *V, uint32_t nV, const int32_t *S, uint32_t nS,
                          snch_scene **out);
/* Batch of n_meshes independent triangle meshes in one scene (no reference
 * counterpart: the reference builds one lbvh::scene<3> per mesh,
 * scene.cuh:1128-1229).  V: vertices of all meshes, concatenated (v_off[m] ..
 * v_off[m+1] - 1 belong to mesh m); F: triangles of all meshes, concatenated
 * (f_off[m] .. f_off[m+1] - 1), with vertex indices LOCAL to the mesh; at
 * most 2^24 triangles per mesh (as for snch_scene_create_tri).
 * v_off and f_off hold n_meshes + 1 entries starting at 0 (host).  Each mesh
 * gets its own edge table, scene box and tree, identical to those of the
 * mesh built alone; one build and one query launch cover all meshes. */
int snch_scene_create_tri_batch(snch_context *ctx, uint32_t n_meshes, const float *V, const uint32_t *v_off,
                                const int32_t *F, const uint32_t *f_off, snch_scene **out);
int snch_scene_destroy(snch_scene *scene);
int snch_scene_mesh_count(const snch_scene *scene, uint32_t *n_meshes);
/* dim (2|3), primitive count, silhouette count (edges in 3D, vertices in 2D) */
int snch_scene_info(const snch_scene *scene, uint32_t *dim, uint32_t *n_prims, uint32_t *n_silhouettes);

/* scene<3>::build_bvh() -> bvh::construct() (scene.cuh:1205-1229, bvh.cuh:380-613),
 * entirely on the GPU, for every mesh of the scene at once.  opts->leaf_size
 * k in 1..8: k = 1 is the reference's tree; k > 1 keeps the same tree and
 * lets the traversal queries treat any internal node over <= k primitives as
 * one leaf (results do not change).  Synchronises `stream` before returning. */
int snch_scene_build(snch_scene *scene, const snch_build_opts *opts, void *stream);
/* Milliseconds of GPU time of the last build, measured with HIP events. */
int snch_scene_build_time(const snch_scene *scene, float *ms);
/* GPU milliseconds of the last build's phases, when it was made with
 * SNCH_BUILD_TIME_PHASES (zeros otherwise): ms[0] leaf boxes, cones and silhouette
 * records, scene box and Morton codes; ms[1] radix sort; ms[2] Karras emission;
 * ms[3] refit (with the gather of the leaves into sorted order).  No counterpart
 * in the reference, which never times its build (benchmark.cu:374-381). */
int snch_scene_build_phases(const snch_scene *scene, float ms[4]);

/* ---- queries (device pointers; asynchronous on `stream`) ---------------- */
/* query_device(bvh, nearest(p), distance_calculator) (query.cuh:238-318).
 * q: n x dim floats.  out_idx: primitive index, out_dist: distance. */
int snch_closest_point(const snch_scene *scene, const float *q, uint32_t n, uint32_t *out_idx, float *out_dist,
                       void *stream);
/* query_device(bvh, nearest_silhouette(p, flip), silhouette_distance_calculator)
 * (query.cuh:325-423).  The reference returns only the distance; out_sil (may
 * be NULL) also receives the silhouette id (edge id in 3D, vertex id in 2D).
 * seed_radius (may be NULL): per-query pruning cap r0; the search is exact
 * within r0 and falls back to the full search when nothing is found there, so
 * outputs equal those of the unseeded call. */
int snch_closest_silhouette(const snch_scene *scene, const float *q, uint32_t n, int flip, const float *seed_radius,
                            uint32_t *out_sil, float *out_dist, void *stream);
/* query_device(bvh, ray_intersect<TestOnly>(ray(o, d), tmax), intersect_test)
 * (query.cuh:79-169).  tmax may be NULL (= +inf).  any_hit selects TestOnly:
 * only out_hit is meaningful.  out_idx/out_t/out_uv may be NULL. */
int snch_ray(const snch_scene *scene, const float *o, const float *d, const float *tmax, uint32_t n, int any_hit,
             uint8_t *out_hit, float *out_t, float *out_uv, uint32_t *out_idx, void *stream);
/* sample_object_in_sphere(bvh, sphere_intersect(c, r), intersect_sphere,
 * measurement_getter, green_weight, u[0]) then, if out_pt != NULL,
 * sample_on_object(bvh, idx, sample_on_object, u[1], u[2]) (sample.cuh:7-92).
 * c: n x dim, r: n, u: n x 3.  out_idx = -1 and out_pdf = 0 when nothing is
 * sampled. */
int snch_sample_in_sphere(const snch_scene *scene, const float *c, const float *r, const float *u, uint32_t n,
                          int32_t *out_idx, float *out_pdf, float *out_pt, void *stream);
/* sample_on_object(bvh, idx, sample_on_object, u, v) (sample.cuh:7-21; scene.cuh:1231-1240) */
int snch_sample_on_object(const snch_scene *scene, const int32_t *idx, const float *uv, uint32_t n, float *out_pt,
                          void *stream);
/* query_device(bvh, overlaps(box), out, max) (query.cuh:176-231; predicator.cuh:115-136):
 * for each query box [lo[i], hi[i]] (n x dim floats each), the primitives whose
 * leaf boxes intersect it (aabb.cuh:54-82), in the reference's visit order.
 * The first max_per_query ids land in out_idx[i * max_per_query ...];
 * out_count[i] is the number found (it may exceed max_per_query, as the
 * reference's return value does).  out_idx may be NULL when max_per_query = 0. */
int snch_overlap(const snch_scene *scene, const float *lo, const float *hi, uint32_t n, uint32_t max_per_query,
                 uint32_t *out_count, uint32_t *out_idx, void *stream);
/* query_device(bvh, line_intersect(line(o, d)), element_test, out, max)
 * (query.cuh:12-77; predicator.cuh:8-40; line-box test aabb.cuh:330-364): the
 * primitives the infinite line o + t d crosses, t of any sign, with the
 * scene's ray test (scene.cuh:1005-1052 / :545-576) as the element test minus
 * its t >= 0 condition; ids and line parameters t in the reference's visit
 * order, as for snch_overlap.  out_t may be NULL. */
int snch_line_intersect(const snch_scene *scene, const float *o, const float *d, uint32_t n, uint32_t max_per_query,
                        uint32_t *out_count, uint32_t *out_idx, float *out_t, void *stream);
/* Same queries with explicit traversal options. */
/* Morton visiting order of n query points (dim floats each) in the scene's
 * root box: out_perm (DEVICE uint32[n]) receives a permutation of [0, n).
 * No reference counterpart: the reference visits queries in input order. */
int snch_query_order(const snch_scene *scene, const float *q, uint32_t n, uint32_t *out_perm, void *stream);
int snch_closest_point_ex(const snch_scene *scene, const float *q, uint32_t n, uint32_t *out_idx, float *out_dist,
                          const snch_query_opts *opts, void *stream);
int snch_ray_ex(const snch_scene *scene, const float *o, const float *d, const float *tmax, uint32_t n, int any_hit,
                uint8_t *out_hit, float *out_t, float *out_uv, uint32_t *out_idx, const snch_query_opts *opts,
                void *stream);
int snch_closest_silhouette_ex(const snch_scene *scene, const float *q, uint32_t n, int flip,
                               const float *seed_radius, uint32_t *out_sil, float *out_dist,
                               const snch_query_opts *opts, void *stream);

/* ---- batched queries over every mesh of a batch (no reference counterpart;
 * the reference launches one kernel per mesh and query set) ---------------
 * q: all query points (DEVICE, n x 3 floats), grouped by mesh: the queries of
 * mesh m are q_off[m] .. q_off[m+1] - 1 (q_off: HOST, n_meshes + 1 entries,
 * q_off[0] = 0).  Outputs are per query, in input order; ids are local to the
 * query's mesh.  One persistent launch: waves take chunks of
 * opts->queries_per_wave queries of one mesh from a work queue (meshes with
 * more primitives first).  Results equal the per-mesh single calls. */
int snch_batch_closest_point(const snch_scene *scene, const float *q, const uint32_t *q_off, uint32_t *out_idx,
                             float *out_dist, const snch_query_opts *opts, void *stream);
/* seed_radius (may be NULL): per-query pruning radius (DEVICE, n floats), as
 * for snch_closest_silhouette: the search within it is exact and a query not
 * settled there searches again without it, so outputs equal the unseeded call's. */
int snch_batch_closest_silhouette(const snch_scene *scene, const float *q, const uint32_t *q_off, int flip,
                                  const float *seed_radius, uint32_t *out_sil, float *out_dist,
                                  const snch_query_opts *opts, void *stream);

/* Autotuner for the batched queries (kind 0 = closest point, 1 = closest
 * silhouette): coordinate descent from leaf 1 / block 64 / stack 16 / 64
 * queries per wave over leaf size {1, 2, 4, 8} (rebuilding), block size {64,
 * 128, 256, 512}, stack depth {8, 16, 24, 32} and queries per wave {64, 128,
 * 256}, each axis swept with the others at their best so far; one warm-up
 * and one HIP-event-timed launch per configuration over all queries
 * (sample_per_mesh = 0) or the first sample_per_mesh queries of each mesh.
 * Leaves the scene built with the best leaf size.  Pass the result's
 * block_size, stack_depth and queries_per_wave in snch_query_opts.
 * Synchronises. */
typedef struct snch_tune_result {
    uint32_t leaf_size, block_size, stack_depth, queries_per_wave;
    float ms;         /* best time of one batched launch over the sample */
    float ms_default; /* the same for leaf 1, block 64, stack 16, 64 queries per wave */
    /* n_tried: configurations timed */
    uint32_t n_tried;
} snch_tune_result;
int snch_batch_autotune(snch_scene *scene, uint32_t kind, const float *q, const uint32_t *q_off,
                        uint32_t sample_per_mesh, int flip, snch_tune_result *out, void *stream);

/* ---- inspection (host outputs) ------------------------------------------ */
/* The tree in the reference's layout (bvh.cuh:27-33, :625-633): nodes are
 * (parent, left, right, object) per node, internal nodes 0..N-2, leaves
 * N-1..2N-2; boxes are (lo xyz, hi xyz); cones (axis xyz, half_angle, radius)
 * in 3D or (axis xy, half_angle, radius) in 2D.  Any output may be NULL. */
int snch_scene_export(const snch_scene *scene, uint32_t *nodes, float *boxes, float *cones, uint32_t *sorted_codes,
                      uint32_t *sorted_idx);
/* The same for mesh `mesh` of a batch (ids local to the mesh). */
int snch_scene_export_mesh(const snch_scene *scene, uint32_t mesh, uint32_t *nodes, float *boxes, float *cones,
                           uint32_t *sorted_codes, uint32_t *sorted_idx);
/* A built mesh's tree for users' own HIP kernels: the reference's
 * get_bvh_device_ptr() (scene.cuh:1242-1252) feeding query_device calls from
 * device code (query.cuh:79-423, sample.cuh:7-92).  Pass it to
 * lbvh::device::bvh_view (include/snch_lbvh/device.hpp).  All pointers are
 * device pointers into the scene, valid until its next build or destruction.
 * For a batch, `mesh` picks the mesh (ids local to it); else pass 0. */
typedef struct snch_device_view {
    uint32_t dim;      /* 2 or 3 */
    uint32_t n_leaves; /* primitives of the mesh */
    uint32_t flags;    /* traversal options recorded by the build */
    uint32_t qbase;    /* stack key quantisation base of the tree */
    const void *tnode, *tcone, *tleaf, *tedge, *nbox, *ncone; /* float4 records */
    const void *V;     /* vertices, float4 */
    const int32_t *F;  /* the mesh's primitives' vertex indices into V */
} snch_device_view;
int snch_scene_device_view(const snch_scene *scene, uint32_t mesh, snch_device_view *out);

/* Edge (3D) or vertex (2D) silhouette table: slots (opp0, lo, hi, opp1) per
 * edge / (prev, v, next, -1) per vertex, and per-primitive owned ids (3 per
 * triangle, 2 per segment, -1 padded).  For a batch: the tables of the meshes
 * concatenated, with mesh-local vertex and edge ids. */
int snch_scene_silhouettes(const snch_scene *scene, int32_t *slots, int32_t *owned);

/* ---- learned distance bound (Fourier-feature MLP on BF16 MFMA) ----------
 * No counterpart in the reference.  3 -> 64 Fourier features (sigma) -> 4 x 128
 * SiLU -> 1, BF16 weights and activations, FP32 accumulation.  It regresses a
 * query's closest-silhouette distance; snch_mlp_predict(mode = 1) turns the
 * prediction into a seed radius for snch_closest_silhouette, whose exact
 * verification and fallback keep results identical to the unseeded call. */
typedef struct snch_mlp snch_mlp;
uint32_t snch_mlp_num_params(void);
/* lo/hi: the box the inputs live in (normalised to [-1, 1]^3) */
int snch_mlp_create(snch_context *ctx, const float *lo, const float *hi, float sigma, uint64_t seed, snch_mlp **out);
int snch_mlp_destroy(snch_mlp *mlp);
/* host copies: params_out/params_in (snch_mlp_num_params() floats, FP32 master
 * weights), fourier_out (B, 32 x 3: features sin(B xhat), cos(B xhat)), frame_out (8 floats:
 * centre xyz, 1/half-extent xyz, 1/D, D).  Any may be NULL. */
int snch_mlp_params(snch_mlp *mlp, float *params_out, const float *params_in, float *fourier_out, float *frame_out);
/* Adam training on device points (n x 3) and exact distances (n); batch must be
 * a multiple of 64; epoch_loss (host, epochs floats, may be NULL) gets the mean
 * loss per epoch.  Synchronises the stream. */
int snch_mlp_train(snch_mlp *mlp, const float *pts, const float *dist, uint32_t n, uint32_t batch, uint32_t epochs,
                   float lr, float *epoch_loss, void *stream);
/* Loss: w (o - t)^2 with t = log(min(d / D, 1) + 1e-3), w = 2 tau when the
 * prediction o is below t and 2 (1 - tau) above it (an expectile; tau = 0.5,
 * the default, is the squared error).  tau > 0.5 trains an upper bound. */
int snch_mlp_set_loss(snch_mlp *mlp, float tau);
/* Gradient of the mean loss over n points (host output, FP32), for checking. */
int snch_mlp_gradient(snch_mlp *mlp, const float *pts, const float *dist, uint32_t n, float *grad_out,
                      float *loss_out, void *stream);
/* mode 0: predicted distance; mode 1: seed radius (pred * (1 + rel) + abs * D) */
int snch_mlp_predict(snch_mlp *mlp, const float *pts, uint32_t n, int mode, float rel, float abs_margin, float *out,
                     void *stream);

/* ---- wavefront walk-on-stars (Laplace, Dirichlet g = x*y + zero Neumann) --
 * No counterpart in the reference: the consumer that drives all four queries
 * per step over the shrinking set of live walks.  The step loop can be
 * captured once as a hipGraph and replayed (use_graph = 1). */
typedef struct snch_wost snch_wost;
/* neumann may be NULL (pure Dirichlet problem); scenes must be built 3D scenes */
int snch_wost_create(snch_context *ctx, const snch_scene *dirichlet, const snch_scene *neumann, uint32_t max_walks,
                     snch_wost **out);
int snch_wost_destroy(snch_wost *wost);
/* x0: device n x 3 start points inside the domain; out: device n estimates;
 * live_after (host, may be NULL): walks still live after max_steps (0). */
int snch_wost_solve(snch_wost *wost, const float *x0, uint32_t n, uint32_t max_steps, float eps, uint64_t seed,
                    int use_graph, float *out, uint32_t *live_after, void *stream);

/* ---- the closest-silhouette query without device allocation -------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, results and visiting order as
 * snch_closest_silhouette (Morton order from SNCH_ORDER_MIN_QUERIES queries
 * up), but the ordering scratch lives in a static device array of each GPU
 * (about 71 MB, sized for SNCH_NOALLOC_MAX_QUERIES queries, held until the
 * process ends) instead of per-scene buffers grown with hipMalloc; n above
 * SNCH_NOALLOC_MAX_QUERIES is rejected with SNCH_ERR_INVALID.  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller.  With a
 * seed radius, queries not settled within it are searched again at once by
 * their own lane (snch_query_opts.seed_retry = 1), which needs no queue. */
#define SNCH_NOALLOC_MAX_QUERIES (1u << 22)
int snch_closest_silhouette_noalloc(const snch_scene *scene, const float *q, uint32_t n, int flip,
                                    const float *seed_radius, uint32_t *out_sil, float *out_dist, void *stream);

/* ---- device memory helpers for callers without their own allocator ------- */
int snch_device_malloc(void **ptr, uint64_t bytes);
int snch_device_free(void *ptr);
int snch_memcpy_h2d(void *dst, const void *src, uint64_t bytes, void *stream);
int snch_memcpy_d2h(void *dst, const void *src, uint64_t bytes, void *stream);
/* Waits for `stream`, then reports (and clears) the device's sticky error
 * words: SNCH_ERR_OVERFLOW if a traversal lost a stack entry, SNCH_ERR_HIP if a
 * radix-sort pass exceeded its wait bound (neither is expected; snch_scene_build
 * checks them too). */
int snch_stream_synchronize(void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SNCH_H */
